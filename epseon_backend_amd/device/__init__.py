"""Interfaces of compute devices."""
from __future__ import annotations
