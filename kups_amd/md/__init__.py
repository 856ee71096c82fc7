"""Molecular dynamics: state construction and fused integrator propagators."""
