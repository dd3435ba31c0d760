"""FP64 CPU oracle of the MGBF localization operator (test infrastructure only).

Only tests/, the smoke check and the CPU-baseline leg of bench.py may import
this package, and only as the checker. The product path never routes through it.
"""
