"""CPU oracle for the SLOS strong-simulation path — TEST INFRASTRUCTURE ONLY.

Nothing in the product package (`photonic_engines_amd`) imports, links or runs anything under
this directory. Only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of
`bench.py` use it, and only as the checker or as the timed CPU baseline.

Contents:
  slos_oracle.py  numpy restatement of the reference algorithm: layer enumeration in the
                  reference order, stored raising edges, the complex128 stored-edge recurrence,
                  the readout, uniform-loss thinning, Ryser permanents. Every function cites the
                  reference file:line it restates.
  pull_oracle.c   C / OpenMP pull-form restatement (complex128 or complex64, no stored tables)
                  and a Gray-code Ryser permanent; built into _build/liboracle.so by
                  oracle/Makefile. Used for medium/large parity checks and as the CPU baseline.

Pinning (what each function is checked against, in tests/test_oracle.py):
  enumerate_layers  brute force over occupation vectors + a recursive builder (independent
                    code), closed-form sizes, itertools order for fock/unbunched.
  edges             brute-force creation-operator rule on occupation vectors.
  slos_*            Ryser permanents (which are themselves pinned by the permutation sum and
                    per(ones(n)) = n!), sum p = 1, U = I, permutation U, Hong-Ou-Mandel.
  thin_distribution closed-form binomial masses, eta = 1 identity, the independent explicit
                    pattern route (thin_patterns restatement), brute-force photon deletion.
  pull_oracle.c     the numpy oracle on every mask and bunched inputs.
"""
