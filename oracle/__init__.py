"""Oracle: CPU restatement of the reference aligners (test infrastructure only).

Only tests/, the smoke check of __graft_entry__ and the cpu_baseline leg of
bench.py import this package, as the checker and as the CPU timing.
"""
