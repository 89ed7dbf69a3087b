"""Solver policy: tolerances and algebraic multigrid."""
