"""Host-side mirror of the reference's bake entry points over the C-ABI library."""
from ._lib import BakeError, Grid, Job  # noqa: F401
