"""MI355X-native exact kNN with GSKNN's dgsknn/sgsknn C API.

The product is the C-ABI library rnn_amd/lib/libgsknn.so (HIP kernels for
gfx950); ``rnn_amd.gsknn`` is its Python binding, mirroring the reference's
python/gsknn.py.
"""
