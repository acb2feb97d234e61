"""CPU oracle for the DMC propagation path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (``pyvibdmc_amd``) may
import, call or link anything in this package.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker or as the timed CPU baseline.

Contents
--------
``dmc``        FP64 NumPy restatement of ``DMC_Sim``'s step loop with injectable
               random draws (reference ``pyvibdmc/pyvibdmc.py``).
``pes``        NumPy potential-energy surfaces and trial wave functions used by
               the configurations, with analytic derivatives.
``mlp``        NumPy float32 restatement of the surrogate MLP (forward, backward,
               Adam), with bfloat16 rounding emulation.
``eigen``      Exact discrete-time-step ground-state energy of the reference's
               branching scheme for a 1-D harmonic oscillator (grid eigenvalue).
``initial``    The reference's element-by-element finite-difference Hessian
               (initial_conditioner/harmonic_analysis.py) and the Wilson GF-matrix
               frequencies of a bent triatomic that pin it.
``cpu_ref``    The CPU baseline: the oracle driven like the reference's
               ``Potential`` manager (PES fanned out over a process pool).

Pinning status of each function is stated in its docstring; every function
listed here is checked in ``tests/`` against golden vectors written by the
reference itself (``tests/golden/``), closed forms, or brute force.
"""
