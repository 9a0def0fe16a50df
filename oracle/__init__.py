"""ORACLE — test infrastructure only.

A CPU restatement of Pippin's FP64 engine (C++, built by oracle/Makefile into
oracle/_build/liborc.so) plus a Python restatement of its MJCF loader. Only
tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may import
this package, and only as the checker. The product (pippin_amd/) never does.
"""
