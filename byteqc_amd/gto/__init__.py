from byteqc_amd.gto.mole import (Mole, M, BOHR, gto_norm, cart2sph, cart_components,
                                 load_basis, parse_nwchem)
from byteqc_amd.gto.mole import (ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF, PTR_EXP,
                                 PTR_COEFF, CHARGE_OF, PTR_COORD, ATM_SLOTS, BAS_SLOTS)
