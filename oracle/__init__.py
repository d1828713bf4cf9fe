"""CPU FP64 oracle of the film-height step. Test infrastructure: only the
tests, the smoke check and the CPU-baseline leg of bench.py load it."""
