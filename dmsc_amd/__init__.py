"""MI355X-native Discrete Morse-Smale complex (drop-in for the reference ``dmsc`` GPU path)."""

from .dmsc import MSComplex as MSComplex
from .dmsc import compute_dmsc as compute_dmsc
from .dmsc import compute_dmsc_batch as compute_dmsc_batch
from .diff import persistence_loss as persistence_loss
from .synthetic import generate_noisy_landscape as generate_noisy_landscape

__all__ = ["compute_dmsc", "compute_dmsc_batch", "generate_noisy_landscape", "MSComplex", "persistence_loss"]
