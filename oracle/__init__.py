"""ORACLE package: CPU restatement of the reference BP, used only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, as the reference the
HIP results are checked against (never as a product path)."""
