"""Python side of the MI355X k-space SBE propagator (libwannier.so).

Modules:
  lib       locate and load the native library; the C-ABI symbol list
  tbgen     W90 _tb.dat writer, Haldane honeycomb and seeded synthetic models
  spectrum  HHG spectrum and per-harmonic intensities from J(t)
"""
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libwannier.so")
