"""CPU restatements of FEMFlow's MLS-MPM substep: the checker for the HIP path.

Test infrastructure only -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the femflow_amd package."""
