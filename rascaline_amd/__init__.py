"""MI355X-native SOAP / LODE descriptor hot path with featomic's calculator API."""
