"""ctypes driver for libwannier.so with the reference's pySWE interface."""
