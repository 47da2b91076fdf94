"""CPU oracle of the wind-tunnel solver path (test infrastructure only)."""
