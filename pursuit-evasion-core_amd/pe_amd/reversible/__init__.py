"""Reversible-car games: the free-space 5D relative HJI (free_dp) and the
dense 8D square game (world_dp, read back by world_table)."""
