"""Core data model of the kups_amd host layer (constants, batches, neighbor lists)."""
