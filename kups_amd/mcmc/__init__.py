"""Grand-canonical Monte Carlo of rigid molecules (GCMC) on the GPU."""
