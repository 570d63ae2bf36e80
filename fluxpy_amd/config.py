"""Package-wide flags (config.py:1-2 of the reference)."""
DEBUG = False
DEFAULT_EPS = 1e-5
