"""TEST INFRASTRUCTURE ONLY: CPU oracles of the reference MindQuantum mqvector path.

``oracle.cpu`` binds the C++ restatement (oracle/mqvector_oracle.cpp, built by oracle/Makefile
into oracle/_build/liboracle.so); ``oracle.dense`` is an independent dense numpy simulation for
small registers.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use them.
"""
