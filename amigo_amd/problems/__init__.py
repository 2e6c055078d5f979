"""Seeded workload generators (component-group NLPs on the device)."""
