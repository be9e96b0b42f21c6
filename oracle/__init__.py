"""Oracle: CPU restatement of the reference's H_eff hot path and its sweep, used only as a checker by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg. Nothing in dmrg_amd imports it."""
