from .g2p import g2p
from .grid_op import check_collision_points, grid_op
from .p2g import p2g
