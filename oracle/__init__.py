"""Test infrastructure only: the CPU restatement of the reference's RHS + RK3
path, used as the checker by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  Nothing in the product imports it."""
