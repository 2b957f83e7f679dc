"""MI355X (gfx950) back end for HORA's zone-graph engine: network IR and host API over the C ABI."""
from .network import Network, Target  # noqa: F401
