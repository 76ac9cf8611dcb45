__version__ = "0.1.11+mi355x"
