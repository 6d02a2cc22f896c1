"""MI355X-native multislice engine with MULTEM's ilc_multem call surface."""
from .parameters import Parameters, SystemConf  # noqa: F401
from . import crystal, configs  # noqa: F401
