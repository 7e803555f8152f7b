"""CPU oracle for the Weft tile-fit scoring path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this package, and only as the checker / the timed CPU baseline.  The product
package ``weft_amd`` never imports it and never falls back to it.

- ``oracle.weft_oracle.c`` (built into ``liboracle.so`` by ``oracle/Makefile``): float32
  restatement of render.render_tile, _tile_objective, batch_mse, _greedy_add and the
  plain tile fit (see the file header for the reference file:line of each function).
- ``oracle.render_np``: NumPy restatement of the same functions, written in the
  reference's own array style, used to cross-check the C restatement on random inputs.
- ``oracle.coracle``: ctypes bindings for ``liboracle.so``.
"""
