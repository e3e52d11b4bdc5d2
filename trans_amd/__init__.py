"""MI355X-native residual-and-update hot path of the TRANS RANS solver (host side)."""
