"""RRSPH_HIP: an MI355X-native WCSPH time step behind the RackotRR/SPH2D step API."""
