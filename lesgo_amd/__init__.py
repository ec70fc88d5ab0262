"""lesgo_amd: MI355X-native (gfx950 HIP) lesgo LES time step behind a C ABI."""
from .case import Case, Config, aligned_farm, seeded_state  # noqa: F401
from ._lib import FIELDS, FIELD_ID, LesgoError  # noqa: F401
