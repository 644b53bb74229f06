"""Extension-kernel layer (the drop-in boundary of tupan.lib)."""
