"""Subpackage for interacting with CPU compute capabilities."""
from __future__ import annotations
