# Version of the MI355X build; the reference interface it mirrors is imate 0.25.2.
__version__ = "0.1.0"
