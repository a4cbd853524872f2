"""Checkers for the tests and the CPU baseline of bench.py (never the product path)."""
