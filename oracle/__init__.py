"""Host-loop oracle of the stable-fluids 3D step (test infrastructure only)."""
