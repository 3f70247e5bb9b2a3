"""Utilities with the reference's gpaw.utilities module layout."""
