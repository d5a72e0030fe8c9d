"""Single-species DGFS 1D3V solver (dgfsStd1D) on one MI355X."""
