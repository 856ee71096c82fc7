"""The YAML applications (kups_md_lj, kups_mcmc_rigid) over the HIP kernels."""
