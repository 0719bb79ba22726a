"""The oracle: a CPU restatement of jNO's P1/P2 FEM hot path.

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker, never by the jno_amd product path.
"""
