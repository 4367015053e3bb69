__version__ = "0.1.19+mi355x.0.1"
