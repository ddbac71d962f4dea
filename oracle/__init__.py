"""CPU oracle — TEST INFRASTRUCTURE ONLY.

FP64 restatements of the reference algorithms on the hot path, used by the
tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the
checker / timed CPU baseline.  Nothing in the product imports this package.

  fd_oracle.c   forward stepping (wave_equation.cpp:39-152, 262-405;
                fault_tolerance.cpp:49-157, 198-257), Ricker, sponge factors
  fd.py         ctypes wrappers of liboracle.so
"""
