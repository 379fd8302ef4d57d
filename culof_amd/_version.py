"""Single source of the package version (the reference's src/culof/_version.py:8)."""

__version__ = "0.2.0"
