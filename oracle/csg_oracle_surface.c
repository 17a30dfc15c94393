void oracle_surface_placeholder(void) {}
