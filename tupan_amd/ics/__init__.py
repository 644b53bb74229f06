"""Initial conditions (tupan/ics)."""
