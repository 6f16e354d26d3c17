"""ctypes loader of the native library (mirrors cuobc/lib/__init__.py:41-42).

The library is built in-tree by ``python -m byteqc_amd.build`` (or
``__graft_entry__.build()``).  There is no fallback: if libbqcjk.so is
missing, loading raises and the J/K path is unavailable.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
c_int_p = ctypes.POINTER(ctypes.c_int)

_SIGNATURES = {
    'GINTinit_basis_prod_cpu': (None, [c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_int,
                                       c_void_p, c_int, c_void_p, c_int, c_void_p]),
    'GINTinit_basis_prod_gpu': (None, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p]),
    'GINTdel_basis_prod': (None, [c_void_p]),
    'BQCdel_basis_prod_clone': (None, [c_void_p]),
    'GINTbuild_jk': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p,
                             c_void_p, c_int, c_int, c_int, c_void_p]),
    'BQCbuild_jk': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                            c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p,
                            c_double, c_void_p]),
    'BQCbuild_jk_noalloc': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                    c_int, c_void_p, c_void_p, c_int, c_void_p, c_int,
                                    c_void_p, c_void_p, c_double, c_void_p]),
    'BQCq_cond': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int,
                          c_void_p, c_void_p]),
    'BQCint1e': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                         c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    'BQCbuild_jk_batch': (c_int, [c_int] + [c_void_p] * 5 + [c_int, c_void_p, c_void_p,
                                  c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p,
                                  c_double, c_void_p]),
    'BQCgraph_capture': (c_void_p, [c_int] + [c_void_p] * 5 + [c_int, c_void_p, c_void_p,
                                    c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p,
                                    c_double, c_void_p]),
    'GINTfill_int2e': (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_int, c_int, c_int, c_void_p]),
    'BQCgraph_launch': (c_int, [c_void_p, c_void_p]),
    'BQCgraph_destroy': (None, [c_void_p]),
    'BQCset_tile': (c_int, [c_int, c_int, c_int, c_int]),
    'BQCget_tile': (c_int, [c_int, c_int, c_int, c_int]),
    'BQCset_launch_mode': (c_int, [c_int]),
    'BQCset_plan_threads': (c_int, [c_int]),
    'BQCset_row_split': (c_int, [c_int, c_int]),
    'BQCplan_tasks': (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_double,
                              c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    'BQCdm_cond': (c_int, [c_void_p] * 7 + [c_int, c_int, c_int, c_void_p]),
    'BQCsymmetrize_jk': (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
    'BQClast_build_flops': (None, [c_void_p]),
    'BQCprofile_enable': (None, [c_int]),
    'BQCprofile_count': (c_int, []),
    'BQCprofile_read': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
}


def library_path(name='libbqcjk'):
    return os.path.join(_HERE, name + '.so')


def load_library(name='libbqcjk'):
    """Load lib/<name>.so with its argument types; raise if it is not built."""
    if name not in _LIBS:
        path = library_path(name)
        if not os.path.exists(path):
            raise RuntimeError(
                f'{path} is not built; run `python -m byteqc_amd.build` first')
        lib = ctypes.CDLL(path)
        for fn, (res, args) in _SIGNATURES.items():
            f = getattr(lib, fn)
            f.restype = res
            f.argtypes = args
        _LIBS[name] = lib
    return _LIBS[name]
