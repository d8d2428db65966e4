"""vapor_amd - MI355X (gfx950) MAC 3D smoke step behind vapor's solver API."""
