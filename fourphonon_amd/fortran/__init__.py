"""Fortran drop-in bridge (module fp4_processes) and the small RTA driver."""
