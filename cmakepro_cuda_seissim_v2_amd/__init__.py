"""MI355X-native 3D/2D elastic velocity-stress solver (Seis3D/Seis2D session API over a HIP C ABI)."""
__version__ = "0.1.0"
