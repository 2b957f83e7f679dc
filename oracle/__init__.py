"""The oracle: a CPU C++ restatement of HORA's zone algebra and forward search.

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import this package; the HIP back end (ptctcl_amd) never does.
"""
