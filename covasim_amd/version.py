__version__ = "0.1.0"  # tracks the behaviour of Covasim 3.1.7
