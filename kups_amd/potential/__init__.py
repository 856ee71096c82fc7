"""Potentials (LJ, Morse, Ewald) with the reference's Potential interface."""
