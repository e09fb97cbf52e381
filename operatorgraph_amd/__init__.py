"""MI355X-native shape-grammar derivation engine (host-side package)."""
