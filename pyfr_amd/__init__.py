"""MI355X-native flux-reconstruction RHS core with PyFR's backend API."""
__version__ = "0.1.0"
