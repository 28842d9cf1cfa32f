"""Python binding of the MI355X DMRG hot-path library (libttgpu.so)."""
from .flat import FlatTensor, Index, make_index, legal_blocks, random_tensor  # noqa: F401
