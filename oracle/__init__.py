"""The oracle: a C restatement of PyFastFlow's flow-routing and GraphFlood
path, used only as a checker (tests/, __graft_entry__.smoke) and as the CPU
baseline of bench.py.  It is never imported by pyfastflow_amd."""
