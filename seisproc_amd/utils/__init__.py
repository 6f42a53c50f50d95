"""Job-path utilities around the migration kernels: the reference's on-disk
output format (migration_output_storage) and its Zarr v2 store (zarr_v2)."""
