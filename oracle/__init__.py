"""TEST INFRASTRUCTURE ONLY: the CPU restatement (oracle) of the reference time step.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package.  The product (planktonindividuals.jl_amd, libpi_hip.so) never does.
"""
