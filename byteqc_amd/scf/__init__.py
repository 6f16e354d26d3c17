from byteqc_amd.scf.hf import get_jk, _VHFOpt, basis_seg_contraction
