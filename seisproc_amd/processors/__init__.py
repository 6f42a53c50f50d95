"""Processors package (migration path only)."""
