"""Kirchhoff PSTM on gfx950 (drop-in for SeisProc's processors/migration)."""
