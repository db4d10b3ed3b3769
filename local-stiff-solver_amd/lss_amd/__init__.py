"""MI355X-native hybrid solver for locally stiff SPD systems (host-side Python front end)."""
