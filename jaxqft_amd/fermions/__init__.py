"""Fermion building blocks (gamma algebra and the native Wilson operator);
same exports as jaxqft/fermions/__init__.py:1-15."""

from .gamma import build_euclidean_gamma, check_gamma_algebra, gamma5
from .wilson import WilsonDiracOperator

__all__ = ["build_euclidean_gamma", "gamma5", "check_gamma_algebra", "WilsonDiracOperator"]
