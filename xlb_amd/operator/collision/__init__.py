from xlb_amd.operator.collision.collision import Collision as Collision
from xlb_amd.operator.collision.collision import BGK as BGK
from xlb_amd.operator.collision.collision import KBC as KBC
from xlb_amd.operator.collision.collision import ForcedCollision as ForcedCollision
from xlb_amd.operator.collision.collision import SmagorinskyLESBGK as SmagorinskyLESBGK
