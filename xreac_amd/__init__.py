"""MI355X-native ReaxFF energy / QEq / fixed-charge force path, behind the xreac API."""

from .calculator import Calculator, Evaluation, Relaxation
from .forcefield import ForceField

__all__ = ["Calculator", "Evaluation", "Relaxation", "ForceField"]
