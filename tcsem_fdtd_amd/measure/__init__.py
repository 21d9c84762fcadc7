"""Measurement tooling: the fixed-step profiling driver used by the rocprof summaries and the HBM
streaming probe that bounds what the volume march can reach.  Not part of the library build."""
