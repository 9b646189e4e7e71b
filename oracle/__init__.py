"""FP64 CPU oracle of the DEM step -- test infrastructure, never product code."""
