"""ORACLE package -- test infrastructure only (see sdl_oracle.h)."""
