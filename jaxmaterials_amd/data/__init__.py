"""Microstructure generation."""

from .microstructure import crossply_lame, crossply_place  # noqa: F401

__all__ = ["crossply_lame", "crossply_place"]
