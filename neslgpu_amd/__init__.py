"""MI355X (gfx950) HIP backend of NESL's CVL vector library."""
