"""CPU oracle for the Wilson-fermion path (test infrastructure only; never
imported by the ``jaxqft_amd`` product code)."""
