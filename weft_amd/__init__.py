"""weft_amd -- MI355X (gfx950) implementation of Weft's encoder tile-fit scoring path.

Hot path: rasterize each candidate primitive over a tile, alpha-composite it onto the
tile's current stack canvas, reduce SSE against the target in NumPy's pairwise float32
order, take the first minimum.  Hand-written HIP kernels (weft_amd/csrc) are loaded
through hipModule by the C ABI in include/weft_hip.h and called from Python with ctypes.

Modules
  gpu_render   drop-in for the reference's src/weft/gpu_render.py (same names/contract)
  fit          plain GPU tile fit (device-resident greedy rounds)
  primitives   Primitive / TileRecord, PRIM serialization (kinds 0-5 + 6/7/8)
  records      60-byte device records, code -> record conversion
  quadtree     adaptive quadtree tile producer, packed TileSet
  rd           lambda, byte sizes, bit estimate
  synth        seeded synthetic images of the benchmark configurations
  build        in-tree build of libweft_hip.so
"""
__version__ = "0.1.0"
