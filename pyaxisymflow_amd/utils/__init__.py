"""Utilities next to the path (VTK image output)."""
