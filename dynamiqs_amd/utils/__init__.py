from .general import dag, dissipator, expect, isket, lindbladian, tensor, todm, tracemm
from .operators import create, destroy, displace, eye, momentum, number, position, sigmam, sigmap, sigmax, sigmay, sigmaz
from .states import coherent, coherent_dm, excited, excited_dm, fock, fock_dm, ground, ground_dm
