"""MI355X-native exchange-correlation quadrature engine for ProteinDF's grid XC path."""
