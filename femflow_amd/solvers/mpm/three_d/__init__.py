from .g2p import g2p
from .grid_op import grid_op
from .p2g import p2g
