"""FP64 CPU oracle for the microLC light-curve + chi2 path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, and only to check results / as the CPU
baseline.  The
product package (microsn_amd) never imports it.
"""
